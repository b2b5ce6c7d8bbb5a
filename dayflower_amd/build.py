"""Build the HIP C-ABI library (libdayflower_hip.so) for gfx950 in-tree.

hipcc compiles every source (host C++ and HIP) with the CDNA4 target; packed
FP32 instructions are disabled for all device code.
"""
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdayflower_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["kernels.hip", "api.cpp", "serve.cpp", "bvh.cpp", "perez_host.cpp", "compiled_scene.cpp", "instances.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DFR_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-result", "-Wno-unused-value", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(shlex.quote(c) for c in cmd))
    return r


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out + err))
        elif verbose and (out or err):
            sys.stderr.write(out + err)
    if failed:
        for cmd, msg in failed:
            sys.stderr.write(msg)
        raise RuntimeError("build failed: " + ", ".join(os.path.basename(c[-3]) for c, _ in failed))
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
