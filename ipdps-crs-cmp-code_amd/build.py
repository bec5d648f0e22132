"""Builds libsemblance.so (gfx950) and the drop-in CLI programs in-tree.

    python ipdps-crs-cmp-code_amd/build.py [--jobs N]

Host C++ is compiled with g++ (OpenMP, no FP contraction), device code with
hipcc for gfx950 only.  Everything lands in ipdps-crs-cmp-code_amd/lib/.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["su_io.cpp", "synth.cpp", "runtime.cpp", "tune.cpp", "serve.cpp"]
DEVICE_SOURCES = ["kernels.hip", "de.hip", "lds.hip", "crs_lds.hip", "probe.hip"]
CLI_SOURCES = {"cmp": "cli_cmp.cpp", "crs": "cli_crs.cpp", "crs-de": "cli_crs_de.cpp"}

INC = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROCM, "include")]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__"] + INC
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fhip-fp32-correctly-rounded-divide-sqrt",
                "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-Wno-unused-command-line-argument"] + INC


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [os.path.join(ROCM, "bin", "hipcc")] + DEVICE_FLAGS + ["-c", path, "-o", out]
    else:
        cmd = ["g++"] + HOST_FLAGS + ["-c", path, "-o", out]
    _run(cmd)
    return out


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in HOST_SOURCES + DEVICE_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    lib = os.path.join(LIB, "libsemblance.so")
    tmp = lib + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-fopenmp", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
          "-Wl,--no-undefined"])
    os.replace(tmp, lib)
    for name, src in CLI_SOURCES.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        _run(["g++"] + HOST_FLAGS + [path, "-o", os.path.join(LIB, name), "-L" + LIB, "-lsemblance",
                                     "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    if verbose:
        print("built", lib)
    return lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build(args.jobs, verbose=True)


if __name__ == "__main__":
    main()
