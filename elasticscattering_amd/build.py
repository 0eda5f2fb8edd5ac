"""Builds the gfx950 shared library of this package in-tree.

    python -m elasticscattering_amd.build          # libes_hip.so
    python -m elasticscattering_amd.build --host   # also the es_run host runner

Every HIP source is compiled with -ffp-contract=off (no fused multiply-add
other than the explicit fma() calls of es_math.h, so device FP64 results equal
the host's) and with the packed-FP32 target feature disabled.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libes_hip.so")
RUNNER = os.path.join(PKG, "es_run")

HIP_SOURCES = ["es_kernels.hip", "es_abi.hip", "es_mlp.hip"]
HOST_SOURCES = ["host/es_config.cpp", "host/es_runner.cpp", "host/es_png.cpp"]
RUNNER_MAIN = "host/es_run_main.cpp"

ARCH = os.environ.get("ES_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([cc, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(os.path.splitext(src)[0]) + ".o")
        objs.append(obj)
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                     "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(4, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # es_run: the command-line runner, linked against the library next to it.
    tmp = RUNNER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, RUNNER_MAIN), "-o", tmp,
          "-L", PKG, "-les_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RUNNER)
    return LIB


def build_oracle():
    """The CPU oracle (test infrastructure) and, where the reference tree exists, its _ref build."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])
    if os.path.isdir(os.environ.get("ES_REFERENCE", "/root/reference")):
        _run(["make", "-s", "-C", odir, "ref", "REF=" + os.environ.get("ES_REFERENCE", "/root/reference")])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true", help="also build the CPU oracle")
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    print(build_library(verbose=args.verbose))
    if args.oracle:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
