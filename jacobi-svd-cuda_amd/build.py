"""Build recipe of libjsvd.so (HIP for gfx950) and the C++ drop-in layer.

Compiles every csrc/*.hip and csrc/*.cpp with hipcc for --offload-arch=gfx950 into
jacobi-svd-cuda_amd/lib/libjsvd.so (in-tree, so it travels with the repository snapshot).

Usage:  python jacobi-svd-cuda_amd/build.py [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libjsvd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code (-packed-fp32-ops).
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def sources() -> list[str]:
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    # host-pass warnings about the packed-fp32 feature are expected; surface anything else
    msgs = [ln for ln in (res.stderr or "").splitlines() if ln.strip() and "packed-fp32-ops" not in ln
            and "is not a recognized feature" not in ln]
    if verbose and msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of libjsvd.so failed")
    shutil.move(tmp, LIB)
    build_cpp_test(verbose)
    build_cli(verbose)
    return LIB


CLI_SRC = os.path.join(PKG, "cli", "main.cpp")
CLI_BIN = os.path.join(LIBDIR, "jacobi-svd")


def build_cli(verbose: bool = False) -> str:
    """The command-line front end (the reference CLI's options and reports), g++ linked to libjsvd.so."""
    tmp = CLI_BIN + ".tmp"
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), CLI_SRC, "-o", tmp,
           "-L", LIBDIR, "-ljsvd", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build of the CLI failed")
    shutil.move(tmp, CLI_BIN)
    return CLI_BIN


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
CPP_TEST_BIN = os.path.join(LIBDIR, "test_dropin")


def build_cpp_test(verbose: bool = False) -> str:
    """The C++ drop-in test driver: plain g++ against include/*.hpp, linked to libjsvd.so."""
    tmp = CPP_TEST_BIN + ".tmp"
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC, "-o", tmp,
           "-L", LIBDIR, "-ljsvd", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build of the C++ drop-in test failed")
    shutil.move(tmp, CPP_TEST_BIN)
    return CPP_TEST_BIN


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))


if __name__ == "__main__":
    main()
