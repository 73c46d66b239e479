"""In-tree build of the native libraries.

    python -m mm_propagation_amd.build          # libmmprop_hip.so (+ libmmprop.so drop-in)

libmmprop_hip.so: the HIP kernels and the C ABI of include/mmprop.h, compiled by
hipcc for gfx950 only. libmmprop.so: the C++ drop-in classes of include/mmprop/
(host C++ over the C ABI). Both land in mm_propagation_amd/lib/.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kernels.hip", "solver.cpp", "setup.cpp", "serving.cpp"]
DROPIN_SOURCES = ["dropin/ComplexArray.cpp", "dropin/ComplexArraysContainer.cpp", "dropin/Sparse3DArray.cpp",
                  "dropin/RungeKutta.cpp", "dropin/MultimodePropagation.cpp", "dropin/MultimodePropagation_GPU.cpp",
                  "dropin/FieldOutput.cpp", "dropin/MultimodePropagationBatch.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result", "-fvisibility=hidden",
    # packed FP32 instructions stay disabled in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) of the library's sources and its public header, in a fixed
    order: compiled into libmmprop_hip.so (mmp_source_hash), so that a library built from other
    sources than the tree's is recognised when it is loaded."""
    h = hashlib.sha256()
    names = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp", ".h")))
    for name in names:
        h.update(name.encode())
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(f.read())
    with open(os.path.join(REPO, "include", "mmprop.h"), "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def _compile_hip(src: str, extra=()) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip_library() -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    stamp = [f'-DMMP_SOURCE_HASH="{source_hash()}"']
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda src: _compile_hip(src, stamp), HIP_SOURCES))
    out = os.path.join(LIBDIR, "libmmprop_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_dropin_library() -> str:
    """C++ drop-in classes (host only), linked against libmmprop_hip.so."""
    srcs = [os.path.join(CSRC, s) for s in DROPIN_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return ""
    out = os.path.join(LIBDIR, "libmmprop.so")
    tmp = out + ".tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-std=c++11", "-O2", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(REPO, "include", "mmprop"),
          "-I" + os.path.join(REPO, "include"), "-o", tmp, *srcs, "-L" + LIBDIR, "-lmmprop_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


CPP_TESTS = ["dropin_propagate", "dropin_selftest", "dropin_batch"]


def build_cpp_tests() -> list:
    """The C++ programs under tests/cpp that exercise the drop-in API (tests/cpp/_build/)."""
    src_dir = os.path.join(REPO, "tests", "cpp")
    out_dir = os.path.join(src_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    outs = []
    for name in CPP_TESTS:
        out = os.path.join(out_dir, name)
        _run([cxx, "-std=c++11", "-O2", "-Wall", "-I" + os.path.join(REPO, "include", "mmprop"),
              os.path.join(src_dir, name + ".cpp"), "-o", out, "-L" + LIBDIR, "-lmmprop", "-lmmprop_hip",
              "-Wl,-rpath,$ORIGIN/../../../mm_propagation_amd/lib"])
        outs.append(out)
    return outs


def build_all(verbose: bool = True) -> None:
    lib = build_hip_library()
    if verbose:
        print("built", lib)
    d = build_dropin_library()
    if verbose and d:
        print("built", d)
    for t in build_cpp_tests():
        if verbose:
            print("built", t)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
