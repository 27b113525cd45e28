"""In-tree build of the fpsample_amd native code.

Produces, next to this file:
  libfpsx.so                       HIP kernels (gfx950) + host runtime + C ABI
  _fpsample<EXT_SUFFIX>            pybind11 module linked against libfpsx.so

Device code is compiled with -ffp-contract=off (the distance is summed exactly
as the reference sums it, with no fused multiply-add), with f32 denormals kept,
and with packed FP32 instructions disabled.

Usage: python fpsample_amd/build.py [-v]
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("FPSX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", CSRC, "-I", INCLUDE,
]

SOURCES_HIP = ["fps_kernels.hip", "fps_batch.hip", "fps_batch_pruned.hip", "fps_pruned.hip", "fps_npdu.hip", "fpsx_api.hip",
               "fpsx_batch.hip", "fpsx_server.cpp", "fpsx_spatial.hip", "fpsx_probe.hip", "fpsx_noalloc.hip"]
SOURCES_HOST = ["kdline.cpp"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is expected.
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")


def build(verbose: bool = False, jobs: int = 4) -> None:
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    for s in SOURCES_HIP:
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in SOURCES_HOST:
        cmds.append([CXX, "-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
                     "-pthread", "-Wall", "-I", CSRC, "-I", INCLUDE, "-c", os.path.join(CSRC, s),
                     "-o", _obj(s)])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    lib = os.path.join(HERE, "libfpsx.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *[_obj(s) for s in SOURCES_HIP + SOURCES_HOST], "-pthread", "-Wl,--no-undefined"], verbose)
    os.replace(tmp, lib)

    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mod = os.path.join(HERE, "_fpsample" + ext)
    py_inc = sysconfig.get_paths()["include"]
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall",
          "-I", INCLUDE, "-I", pybind11.get_include(), "-I", py_inc,
          os.path.join(CSRC, "pymodule.cpp"), "-o", mod + ".tmp",
          "-L", HERE, "-lfpsx", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(mod + ".tmp", mod)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(HERE, "libfpsx.so"))
