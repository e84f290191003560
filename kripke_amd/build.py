"""Builds libkripke_amd.so in-tree: HIP kernels for gfx950 (hipcc) plus the
C++ host mirror (g++), linked into one shared library next to this file."""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkripke_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # no implicit contraction: the kernels place their fused multiply-adds explicitly
    "-ffp-contract=off",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result", "-I/opt/rocm/include"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([HIPCC, "-c"] + HIP_FLAGS + ["-o", obj, src])
    for src in sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp"))):
        obj = os.path.join(BUILD, "host_" + os.path.basename(src) + ".o")
        jobs.append([CXX, "-c"] + CXX_FLAGS + ["-o", obj, src])
    objs = [j[j.index("-o") + 1] for j in jobs]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, LIB)
    # the kripke command, linked against the library next to it
    drv = os.path.join(HERE, "kripke")
    _run([CXX] + [f for f in CXX_FLAGS if f != "-fPIC"] + ["-o", drv + ".tmp", os.path.join(CSRC, "driver", "kripke.cpp"),
          "-L" + HERE, "-lkripke_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
