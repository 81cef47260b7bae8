"""Builds libeojoin.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: python -m eating_one_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libeojoin.so")
CLI = os.path.join(PKG, "eo_join")

SOURCES = [
    "scan.hip",
    "partition.hip",
    "hashjoin.hip",
    "nlj.hip",
    "index.hip",
    "gen.hip",
    "util.hip",
    "batch.hip",
    "api.cpp",
    "api_index.cpp",
    "api_batch.cpp",
    "api_io.cpp",
    "plan.cpp",
    "tuner.cpp",
    "tune_table.cpp",
    "status.cpp",
    "serve.cpp",
    "serve_queue.cpp",
    "noalloc.hip",
    "stream.cpp",
]

ARCH = os.environ.get("EO_OFFLOAD_ARCH", "gfx950")

# Files whose text the source hash covers: every C/C++/HIP source and header
# of the library and the public headers.
_HASHED_EXT = (".hip", ".cpp", ".hpp", ".c", ".h")


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) over the library's sources and headers.

    The build embeds it in libeojoin.so (eo_build_source_hash), so a library
    built from other sources than the tree's is detected before it is used.
    """
    import hashlib

    h = hashlib.sha256()
    files = []
    for d in (CSRC, INCLUDE):
        for name in sorted(os.listdir(d)):
            if name.endswith(_HASHED_EXT):
                files.append(os.path.join(d, name))
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        h.update(b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()[:16]

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE,
    "-I", CSRC,
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build libeojoin")
    return path


def _compile(src: str, defines: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), "-x", "hip", *HIPCC_FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compiles every source and links eating_one_amd/libeojoin.so. Returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    missing = [s for s in SOURCES if not os.path.exists(os.path.join(CSRC, s))]
    if missing:
        raise RuntimeError(f"sources listed in SOURCES are missing: {missing}")
    defines = [f'-DEO_SOURCE_HASH="{source_hash()}"']
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, defines), SOURCES))
    tmp = LIB + ".tmp"
    # every symbol must resolve at link time, not at dlopen or first call
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line driver (plain C over the C ABI)
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc and os.path.exists(os.path.join(CSRC, "eo_join_cli.c")):
        cmd = [cc, "-O2", "-Wall", "-I", INCLUDE, os.path.join(CSRC, "eo_join_cli.c"), "-o", CLI + ".tmp",
               "-L", os.path.dirname(LIB), "-leojoin", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"cli build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print(f"built {LIB} from {len(objs)} objects")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
