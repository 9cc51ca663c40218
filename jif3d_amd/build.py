"""Compile libjif3d_amd.so (HIP, gfx950) in-tree.

    python -m jif3d_amd.build            # build
    python -m jif3d_amd.build --asm      # also write device assembly under _build/asm

The library is one shared object with the C ABI of include/jif3d_gravmag.h.
Object files go to _build/ at the repository root; the .so lands beside this file.
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libjif3d_amd.so")
ARCH = "gfx950"

SOURCES = ["api.cpp", "profiler.cpp", "fp64_path.hip", "mixed_path.hip", "storedj.hip",
           "mixed_check.hip", "maggrad_mixed.hip"]

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libjif3d_amd.so")


def _compile(src, asm=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
           "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function", "-c",
           os.path.join(CSRC, src), "-o", obj] + DEVICE_FLAGS
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{res.stderr}")
    out = [obj, res.stderr]
    if asm and src.endswith(".hip"):
        adir = os.path.join(BUILD, "asm")
        os.makedirs(adir, exist_ok=True)
        acmd = [hipcc(), "-x", "hip", "-O3", "-std=c++17", f"--offload-arch={ARCH}",
                "--cuda-device-only", "-S", "-I", INCLUDE, "-I", CSRC,
                os.path.join(CSRC, src), "-o",
                os.path.join(adir, os.path.splitext(src)[0] + ".s")] + DEVICE_FLAGS
        subprocess.run(acmd, check=True, capture_output=True)
    return out


def build(verbose=False, asm=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, asm), SOURCES))
    objs = [r[0] for r in results]
    if verbose:
        for r in results:
            if r[1].strip():
                print(r[1], file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"linking libjif3d_amd.so failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVERS = ["adapter_driver", "bridge_driver"]


def build_driver():
    """The C++ host drivers (the adapter, and the jif3D imp behind the bridge classes),
    linked against libjif3d_amd.so."""
    cxx = shutil.which("g++") or "g++"
    for name in DRIVERS:
        exe = os.path.join(HERE, "host", name)
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE, "-I",
               os.path.join(HERE, "host"), os.path.join(HERE, "host", name + ".cpp"), "-o",
               exe + ".tmp", "-L", HERE, "-ljif3d_amd", "-Wl,-rpath,$ORIGIN/.."]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"building the C++ driver {name} failed:\n{res.stderr}")
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--asm", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(verbose=a.v, asm=a.asm))
