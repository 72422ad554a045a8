"""Builds the HIP library c2a_amd/lib/libc2a_amd.so for gfx950.

Flags that matter for results:
  -ffp-contract=off   every a*b+c rounds twice, on host and device alike, so
                      the GPU reproduces the host-side arithmetic bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
"""
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libc2a_amd.so")
OBJDIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["ccd_dispatch.cpp", "capi.cpp", "scorer.hip"]
CXX_SOURCES = ["bvh.cpp", "compat.cpp"]
# ccd_kernel.hip is compiled once per kernel shape listed in ccd_shapes.h
SHAPE_SOURCE = "ccd_kernel.hip"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def kernel_shapes():
    """(K, CAP, NW, LRN) of every shape in ccd_shapes.h's C2A_SHAPES list."""
    text = open(os.path.join(CSRC, "ccd_shapes.h")).read()
    body = text[text.index("#define C2A_SHAPES(X)"):]
    body = body[:body.index("\n\n")]
    return [tuple(int(v) for v in m)
            for m in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", body)]


def build(verbose=True, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    jobs = jobs or max(1, min(16, os.cpu_count() or 4))
    procs, objs = [], []
    for k, cap, nw, lrn in kernel_shapes():
        obj = os.path.join(OBJDIR, "ccd_kernel_k%d_cap%d_nw%d_lrn%d.o" % (k, cap, nw, lrn))
        cmd = [cc, "-x", "hip", "-c", os.path.join(CSRC, SHAPE_SOURCE), "-o", obj,
               "-DC2A_SK=%d" % k, "-DC2A_SCAP=%d" % cap, "-DC2A_SNW=%d" % nw,
               "-DC2A_SLRN=%d" % lrn] + COMMON + HIP_FLAGS
        objs.append(obj)
        procs.append(cmd)
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        cmd = [cc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + HIP_FLAGS
        objs.append(obj)
        procs.append(cmd)
    for src in CXX_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        cmd = [cc, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON
        objs.append(obj)
        procs.append(cmd)
    running = []
    for cmd in procs:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        if len(running) >= jobs:
            if running.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in running:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
