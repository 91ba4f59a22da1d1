"""Builds kiyosi_amd/libkiyosi_amd.so for gfx950 with hipcc.

Every HIP source under kiyosi_amd/csrc is compiled separately (in parallel)
and linked into one shared library; the Joe-Kuo direction numbers are
embedded from kiyosi_amd/data. Run ``python -m kiyosi_amd.build``.
"""

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libkiyosi_amd.so"
DATA = PKG / "data" / "joe_kuo_6_21201.u32"
ARCH = os.environ.get("KG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # Packed FP32 instructions stay off: hipcc emits none by itself then.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(PKG.parent / "include"),
]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return proc.stderr


def _compile_hip(src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def _compile_table() -> pathlib.Path:
    obj = BUILD / "joe_kuo_table.o"
    _run(["g++", "-O2", "-fPIC", f"-DKG_JOE_KUO_PATH={DATA}", "-c",
          str(CSRC / "joe_kuo_table.cpp"), "-o", str(obj)])
    return obj


def build(jobs: int = 8) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        futures = [pool.submit(_compile_hip, s) for s in sources]
        futures.append(pool.submit(_compile_table))
        objects = [f.result() for f in futures]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objects)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
