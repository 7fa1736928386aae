"""Build liballium_amd.so (C++ host + HIP kernels for gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "liballium_amd.so")
SHIM = os.path.join(HERE, "lib", "liballium_operator.so")
SOURCES = ["csrc/discretization.cpp", "csrc/sweep_api.cpp", "csrc/sweep_kernels.hip", "csrc/box_sweep.hip", "csrc/cnn.hip", "csrc/trt.hip", "csrc/coop_sweep.hip", "csrc/dsa.hip", "csrc/dsa_api.cpp", "csrc/krylov.hip", "csrc/krylov_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value"]


def build(verbose=False, jobs=8):
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    objdir = os.path.join(HERE, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC, *COMMON, "-c", os.path.join(HERE, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stderr=subprocess.PIPE, stdout=subprocess.PIPE), src))
    err = None
    for p, src in procs:
        out, e = p.communicate()
        if p.returncode != 0:
            err = f"hipcc failed on {src}:\n{e.decode()[-4000:]}"
        elif verbose and e:
            sys.stderr.write(e.decode()[-2000:])
    if err:
        raise RuntimeError(err)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    build_operator_shim(verbose)
    return LIB


def build_operator_shim(verbose=False):
    """lib/liballium_operator.so: the C++ adapter (include/allium_operator.hpp)
    behind one C entry, host code built with g++ against liballium_amd.so."""
    libdir = os.path.join(HERE, "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "csrc", "operator_shim.cpp"), "-L" + libdir, "-lallium_amd", "-Wl,-rpath,$ORIGIN",
           "-o", SHIM]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return SHIM


if __name__ == "__main__":
    print(build(verbose=True))
