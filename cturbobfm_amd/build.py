"""Builds the native libraries in-tree (no JIT cache, so the .so travels with the tree).

  cturbobfm_amd/libtbfm_gpu.so   HIP kernels for gfx950 + C-ABI (include/tbfm_gpu.h)
  cturbobfm_amd/turbobfm         `turbobfm input.ini` command-line driver
  oracle/liboracle.so            CPU restatement, the test oracle (tests, bench CPU baseline)
  oracle/_ref/libtbfm_ref.so     reference translation units, only where /root/reference exists
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtbfm_gpu.so")
CLI = os.path.join(HERE, "turbobfm")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("TBFM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no a*b+c fusion, host or device (bit-level parity with the reference
# build); packed-fp32-ops disabled on the device.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stdout.write(r.stdout)
        raise RuntimeError("build step failed: %s" % " ".join(cmd[:3]))
    return r.stdout


def build_gpu(verbose: bool = False) -> None:
    srcs = [os.path.join(CSRC, "tbfm_gpu.hip"), os.path.join(CSRC, "tbfm_host.cpp"),
            os.path.join(CSRC, "tbfm_io.cpp")]
    objs = []
    for s in srcs:
        o = os.path.join(CSRC, os.path.basename(s) + ".o")
        extra = ["-x", "hip"] if s.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + extra + ["-c", s, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "turbobfm_main.cpp"), "-o", CLI,
                                "-L" + HERE, "-ltbfm_gpu", "-Wl,-rpath,$ORIGIN"])
    for o in objs:
        os.remove(o)


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])


def build_all() -> None:
    build_oracle()
    build_gpu()


if __name__ == "__main__":
    build_all()
