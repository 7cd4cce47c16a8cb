"""Builds the package's native parts in place (run on any host with ROCm; no GPU needed to compile):

  lib/libnslab_hip.so   HIP for gfx950: the solver kernels and the C ABI of include/nslab_hip.h
  lib/nslab_napi.node   N-API addon that binds the C ABI for the Node host module js/nslab.js

Usage: python build-me-anything.github.io_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libnslab_hip.so")
NAPI = os.path.join(LIB_DIR, "nslab_napi.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every product and sum rounds separately, as in the reference's JavaScript.
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_lib():
    os.makedirs(LIB_DIR, exist_ok=True)
    src = os.path.join(HERE, "csrc", "nslab_hip.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_napi():
    node_inc = "/usr/include/node"
    if not os.path.exists(os.path.join(node_inc, "node_api.h")):
        return None
    src = os.path.join(HERE, "napi", "nslab_napi.c")
    tmp = NAPI + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I", node_inc, "-I", os.path.join(ROOT, "include"), src,
          "-L", LIB_DIR, "-lnslab_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, NAPI)
    return NAPI


def build():
    build_lib()
    if os.path.exists(os.path.join(HERE, "napi", "nslab_napi.c")):
        build_napi()


if __name__ == "__main__":
    build()
    print("built", LIB, NAPI if os.path.exists(NAPI) else "")
