"""Builds liblt_amd.so (host C++ + gfx950 HIP kernels) in-tree.

Host sources are compiled with g++ against the HIP runtime headers, device
sources with hipcc for gfx950 only; everything is linked with hipcc into
libtensor_amd/_lib/liblt_amd.so.  Run as ``python -m libtensor_amd.build``.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "liblt_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SRC = ["host/lt_core.cpp", "host/lt_sym.cpp", "host/lt_sched.cpp",
            "host/lt_runtime.cpp", "host/lt_contract.cpp", "host/lt_dotprod.cpp",
            "host/lt_blockops.cpp", "host/lt_symmetrize.cpp",
            "host/lt_server.cpp", "host/lt_capi.cpp"]
DEV_SRC = ["device/lt_kernels.hip", "device/lt_ozaki.hip", "device/lt_scratch.hip"]

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-variable",
              "-Wno-unused-but-set-variable", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include")]
# Packed FP32 instructions are disabled for all device code.
DEV_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-D__HIP_PLATFORM_AMD__"]


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _check_no_scratch(src, out):
    """Kernels must not spill to scratch memory (a register-indexed
    descriptor or a spill turns into per-lane memory traffic)."""
    name = None
    bad = []
    for line in out.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split("[")[0].strip()
        elif "ScratchSize [bytes/lane]:" in line:
            n = int(line.split("ScratchSize [bytes/lane]:")[1].split("[")[0].strip())
            if n:
                bad.append("%s (%d bytes/lane)" % (name, n))
    if bad:
        raise RuntimeError("%s: kernels use scratch memory: %s" % (src, ", ".join(bad)))


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + DEV_FLAGS + ["-Rpass-analysis=kernel-resource-usage",
                                     "-c", path, "-o", obj]
        _check_no_scratch(src, _run(cmd))
    else:
        cmd = ["g++"] + HOST_FLAGS + ["-c", path, "-o", obj]
        _run(cmd)
    return obj


def build(verbose=False):
    """Compiles every source (always) and links the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in HOST_SRC + DEV_SRC if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = lib_path() + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread"])
    os.replace(tmp, lib_path())
    if verbose:
        print("built", lib_path())
    return lib_path()


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
