"""Builds libBaryTree_gpu.so in-tree for gfx950.

Host C++ (trees, lists, orchestration, C ABI) is compiled with g++ and
-ffp-contract=off, so the tree boxes and Chebyshev node coordinates are the
reference's unfused IEEE results bit for bit; the kernels are compiled with
hipcc for gfx950.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libBaryTree_gpu.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("BARYTREE_ARCH", "gfx950")

HOST_SRCS = ["bt_host.cpp", "bt_devmem.cpp", "bt_devtree.cpp", "bt_solver.cpp", "bt_vardeg.cpp", "bt_hermite_plan.cpp", "bt_model.cpp",
             "bt_capi.cpp"]
HIP_SRCS = ["bt_tree.hip", "bt_kernels.hip", "bt_learn.hip", "bt_hermite.hip", "bt_degree.hip", "bt_noalloc.hip"]
# measurement tooling for the benchmark, in a library of its own
PROBE_SRC = "bt_probe.hip"
PROBE_LIB = os.path.join(HERE, "libbt_probe.so")

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-Wall", "-Wno-unused-function",
              "-fopenmp"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             # packed FP32 instructions stay off (some of their forms have given
             # wrong results on gfx950 without any error)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # matrix-core results straight into VGPRs (no AGPR copies per pair)
             "-mllvm", "-amdgpu-mfma-vgpr-form",
             "-munsafe-fp-atomics"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s + ".o")
        run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s + ".o")
        run([f"{ROCM}/bin/hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    # g++ links: the host objects use GCC's OpenMP runtime (libgomp)
    run(["g++", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64", "-fopenmp",
         "-Wl,-rpath," + f"{ROCM}/lib"])
    os.replace(tmp, LIB)
    tmp = PROBE_LIB + ".tmp"
    run([f"{ROCM}/bin/hipcc", *HIP_FLAGS, "-shared", os.path.join(CSRC, PROBE_SRC), "-o", tmp])
    os.replace(tmp, PROBE_LIB)
    if verbose:
        print("built", LIB, "and", PROBE_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
