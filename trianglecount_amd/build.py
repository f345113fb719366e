"""Build liblstc.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

    python -m trianglecount_amd.build          # compile every source, link the library

Each ``csrc/*.hip`` is compiled to an object under ``trianglecount_amd/_build``
in parallel, then linked into ``trianglecount_amd/_lib/liblstc.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "liblstc.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    if r.stderr.strip():
        # hipcc's host pass reports the packed-fp32 feature flag as unknown: harmless
        noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if noise:
            sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    shutil.move(tmp, LIB)
    build_cli(jobs)
    return LIB


CLI = os.path.join(PKG, "cli")
BIN = os.path.join(PKG, "_bin")


def _build_tool(src: str) -> str:
    name = os.path.splitext(os.path.basename(src))[0]
    out = os.path.join(BIN, name)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{CLI}", src, "-o", out + ".tmp",
           f"-L{LIBDIR}", "-llstc", "-Wl,-rpath,$ORIGIN/../_lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    shutil.move(out + ".tmp", out)
    return out


def build_cli(jobs: int | None = None) -> list:
    """Host tools mirroring the reference executables (lstc_tc, lstc_part_tc, lstc_halo, lstc_preprocess)."""
    os.makedirs(BIN, exist_ok=True)
    srcs = sorted(os.path.join(CLI, f) for f in os.listdir(CLI) if f.endswith(".cpp"))
    with cf.ThreadPoolExecutor(jobs or min(len(srcs), os.cpu_count() or 4)) as ex:
        return list(ex.map(_build_tool, srcs))


if __name__ == "__main__":
    print(build())
