"""Build recipe of the MI355X library (hipcc for gfx950, g++ for the host side).

    python eu3d-gpu_amd/build.py

produces, in this directory:
  libeu3d.so   -- HIP kernels + the C ABI (include/eu3d.h, include/eu3d_host.h,
                  include/eu3d_surrogate.h)
  host/eu3d_run -- the command-line driver
Objects go to _build/ (not committed).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(REPO, "_build")
LIB = os.path.join(HERE, "libeu3d.so")
RUN = os.path.join(HERE, "host", "eu3d_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EU3D_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the plain FP64 path keeps the reference's no-contraction
# regime (its Makefile:6-7); packed FP32 ops are disabled for all HIP code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
CXX_FLAGS = ["-std=c++17", "-O3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-Wall"]

HIP_SOURCES = ["csrc/eu3d_hip.hip", "csrc/eu3d_mlp.hip"]
HOST_SOURCES = ["host/eu3d_host.cpp", "host/host_abi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        out = _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = RUN + ".tmp"
    _run(["g++"] + CXX_FLAGS + [os.path.join(HERE, "host", "main.cpp"), "-o", tmp, "-L" + HERE, "-leu3d",
                                "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, RUN)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
