"""Builds coffe_amd/libcoffe_engine.so for gfx950 (hipcc for the kernels, gcc for
the host C) in-tree, so that the library travels with the repository."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcoffe_engine.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("COFFE_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["fftlog_kernels.hip", "mu_kernels.hip", "quad_kernels.hip", "cov_kernels.hip", "noalloc.hip",
               "anyfft.hip", "lensing_kernels.hip", "integrated_kernels.hip"]
C_SOURCES = ["engine.c", "background.c", "interp.c", "covariance.c", "refabi.c"]

INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}",
    # keep packed-FP32 instructions out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
CC_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
            f"-I{ROCM}/include", f"-I{INCLUDE}", "-fno-fast-math"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        out = _run([f"{ROCM}/bin/hipcc", *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out:
            print(out)
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        out = _run(["gcc", *CC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([f"{ROCM}/bin/hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-L{ROCM}/lib", "-lamdhip64", "-lpthread", "-lm", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
