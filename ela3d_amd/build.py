"""Compile the gfx950 library (libela3d.so) and the C++ command-line driver in-tree.

Run as ``python -m ela3d_amd.build`` or through ``__graft_entry__.build()``.
Objects go to ``ela3d_amd/_build/``; the shared library lands next to this file
so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libela3d.so")
CLI = os.path.join(HERE, "bin", "ela3d_run")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("ELA3D_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 (v_pk_*_f32) is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}"]

LIB_SOURCES = ["core.cpp", "models.cpp", "capi.cpp", "shot.cpp", "kernels.hip", "tail.hip",
               "serve.cpp", "tune.cpp", "formats.cpp", "noalloc.hip", "fused.hip"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


# Sanitized build of the host code (AddressSanitizer + UndefinedBehaviorSanitizer),
# for running the CPU tests under the sanitizers: the flags apply to the host
# compilation only (device code is built as usual), and the library goes to
# _san/ so it never replaces the in-tree one.  Load it with ELA3D_LIB and preload
# the shared sanitizer runtime (SAN_RUNTIME) into the Python process.
SAN_OBJ = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(HERE, "_san", "libela3d.so")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
SAN_LINK = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-shared-libsan"]


def san_runtime() -> str:
    """The shared ASan runtime of the ROCm clang (it carries the UBSan handlers too)."""
    out = _run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"]).strip().splitlines()
    path = out[-1] if out else ""
    if not os.path.isabs(path):
        res = _run(["/opt/rocm/llvm/bin/clang", "-print-resource-dir"]).strip().splitlines()[-1]
        path = os.path.join(res, "lib", "linux", "libclang_rt.asan-x86_64.so")
    return path


def _compile(src: str, sanitize: bool = False) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(SAN_OBJ if sanitize else OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, *(SAN_HOST if sanitize else []),
           "-x", "hip", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = True, sanitize: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib = SAN_LIB if sanitize else LIB
    os.makedirs(SAN_OBJ if sanitize else OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    sources = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda s: _compile(s, sanitize), sources))
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *(SAN_LINK if sanitize else []), "-o", tmp, *objs])
    os.replace(tmp, lib)
    cli_src = os.path.join(CSRC, "ela3d_run.cpp")
    if os.path.exists(cli_src) and not sanitize:
        _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", cli_src, "-o", CLI, f"-L{HERE}", "-lela3d",
              f"-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {lib}")
    return lib


def build_oracle(verbose: bool = True, sanitize: bool = False) -> str:
    root = os.path.dirname(HERE)
    odir = os.path.join(root, "oracle")
    make = shutil.which("make")
    if sanitize:  # oracle/_san/libela3d_oracle.so, for ELA3D_ORACLE_LIB
        _run([make or "make", "-s", "-B", "-C", odir, "san"])
        out = os.path.join(odir, "_san", "libela3d_oracle.so")
        if verbose:
            print(f"built {out}")
        return out
    if make:
        _run([make, "-s", "-B", "-C", odir])
    else:
        _run(["gcc", "-O3", "-std=gnu99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o",
              os.path.join(odir, "libela3d_oracle.so"), os.path.join(odir, "ela3d_oracle.c"), "-lm"])
    if verbose:
        print(f"built {os.path.join(odir, 'libela3d_oracle.so')}")
    return os.path.join(odir, "libela3d_oracle.so")


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        # python -m ela3d_amd.build --sanitize, then run the CPU tests with the
        # environment it prints
        lib, olib, rt = build(sanitize=True), build_oracle(sanitize=True), san_runtime()
        print(f"LD_PRELOAD={rt} ASAN_OPTIONS=detect_leaks=0 ELA3D_LIB={lib} ELA3D_ORACLE_LIB={olib}")
    else:
        build()
        if "--oracle" in sys.argv:
            build_oracle()
