"""In-tree build of the gfx950 library and its pybind11 module.

    python libffrs_amd/build.py          # builds libffrs_gpu.so and libffrs<ext>.so

Outputs (git-ignored, shipped with the tree):
  libffrs_amd/libffrs_gpu.so                     C ABI of include/ffrs_gpu.h, HIP kernels for gfx950
  libffrs_amd/libffrs.cpython-*.so               pybind11 module `libffrs_amd.libffrs`
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["ctx.hip", "rs256.hip", "rsi16.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FFRS_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are kept out of every device code object.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

GPU_LIB = os.path.join(HERE, "libffrs_gpu.so")
EXT = sysconfig.get_config_var("EXT_SUFFIX")
PY_MODULE = os.path.join(HERE, "libffrs" + EXT)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_gpu_lib():
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + out)
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-Wl,-soname,libffrs_gpu.so", *objs, "-o", tmp])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_py_module():
    import pybind11
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    tmp = PY_MODULE + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", *inc,
          os.path.join(CSRC, "pyffrs.cpp"), f"-L{HERE}", "-l:libffrs_gpu.so", "-Wl,-rpath,$ORIGIN",
          "-o", tmp])
    os.replace(tmp, PY_MODULE)
    return PY_MODULE


def build():
    build_gpu_lib()
    build_py_module()


def is_built():
    return os.path.exists(GPU_LIB) and os.path.exists(PY_MODULE)


if __name__ == "__main__":
    build()
    print("built", GPU_LIB, PY_MODULE)
    sys.exit(0)
